"""Writes the .odb fixtures tests/golden/odb_<deck>_l<step>.odb.gz with the reference's own writer.

For each case the golden particle state of that step (written by the reference solver, see
make_golden.py) is handed to oracle/_ref/ref_odb, which runs the reference's output_total
(CPP_14-2-output.cpp:15-135) on it; the table it writes is kept, gzip-compressed, as the expected
output of the library's .odb writer for the same state. Runs only where that harness is built
(`make -C oracle ref`).

    python tests/golden/make_odb.py
"""
from __future__ import annotations

import gzip
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from fixtures import load_golden  # noqa: E402

REPO = os.path.dirname(os.path.dirname(HERE))
REF_ODB = os.path.join(REPO, "oracle", "_ref", "ref_odb")

# (golden deck, step): a 2D and a 3D soil-structure deck (boundary, soil and structure rows)
CASES = (("ssi_2d_t1", 6), ("ssi_3d_t1", 5))
V3 = ("xp", "ux", "ax", "vx", "interf", "interfss", "strep")
V6 = ("sig", "eps", "adps")
S1 = ("pre", "prea", "prew", "vsig", "veps", "meps", "divq", "divr", "satu", "suct")


def odb_time(dt: float, l: int) -> float:
    """The simulated time at output step l as the reference's loop accumulates it (t += dt)."""
    t = 0.0
    for _ in range(l):
        t = t + dt
    return t


def fixture_path(name: str, l: int) -> str:
    return os.path.join(HERE, f"odb_{name}_l{l}.odb.gz")


SYNTH = ("odb_synthetic", 12345)  # (fixture name, step label)


def synthetic_state(n: int = 12, seed: int = 7):
    """Every output field distinct and nonzero, values spanning many decades and both signs, and
    every particle type, so that each column and the boundary-row substitution is pinned."""
    rng = np.random.default_rng(seed)
    st = {}
    for k, nc in [(k, 3) for k in V3] + [(k, 6) for k in V6] + [(k, 1) for k in S1]:
        mag = 10.0 ** rng.uniform(-12, 9, size=(n, nc))
        st[k] = (mag * rng.choice([-1.0, 1.0], size=(n, nc))).reshape(n, nc) if nc > 1 else (
            mag * rng.choice([-1.0, 1.0], size=(n, 1))).reshape(n)
    st["type"] = np.array([0, 1, 2, 3, 0, 4, 7, 2, 1, 0, 5, 2][:n], dtype=np.int32)
    st["matype"] = rng.integers(0, 60, size=n).astype(np.int32)
    return st


def write_input(path: str, state, l: int, t: float) -> None:
    n = int(np.asarray(state["type"]).shape[0])
    cols = [np.asarray(state[k], dtype=np.float64).reshape(n, 3) for k in V3]
    cols += [np.asarray(state[k], dtype=np.float64).reshape(n, 6) for k in V6]
    cols += [np.asarray(state[k], dtype=np.float64).reshape(n, 1) for k in S1]
    vals = np.ascontiguousarray(np.concatenate(cols, axis=1))
    assert vals.shape == (n, 49)
    ints = np.stack([np.asarray(state["type"]), np.asarray(state["matype"])], axis=1).astype("<i4")
    with open(path, "wb") as f:
        f.write(b"ODBI" + struct.pack("<iid", n, l, t))
        for i in range(n):
            f.write(vals[i].astype("<f8").tobytes())
            f.write(ints[i].tobytes())


def main() -> None:
    if not os.path.exists(REF_ODB):
        sys.exit("oracle/_ref/ref_odb is not built (make -C oracle ref)")
    for name, l in CASES:
        deck, states = load_golden(name)
        t = odb_time(deck.dt, l)
        with tempfile.TemporaryDirectory() as tmp:
            inp = os.path.join(tmp, "state.bin")
            write_input(inp, states[l], l, t)
            subprocess.run([REF_ODB, inp, tmp], check=True)
            with open(os.path.join(tmp, f"data-{l:010d}.odb"), "rb") as f:
                data = f.read()
        with gzip.GzipFile(fixture_path(name, l), "wb", mtime=0) as g:
            g.write(data)
        print(name, l, len(data), "bytes")
    name, l = SYNTH
    st = synthetic_state()
    np.savez(os.path.join(HERE, f"{name}_state.npz"), **st)  # the input, beside the expected output
    with tempfile.TemporaryDirectory() as tmp:
        inp = os.path.join(tmp, "state.bin")
        write_input(inp, st, l, 0.125)
        subprocess.run([REF_ODB, inp, tmp], check=True)
        with open(os.path.join(tmp, f"data-{l:010d}.odb"), "rb") as f:
            data = f.read()
    with gzip.GzipFile(os.path.join(HERE, f"{name}.odb.gz"), "wb", mtime=0) as g:
        g.write(data)
    print(name, l, len(data), "bytes")


if __name__ == "__main__":
    main()

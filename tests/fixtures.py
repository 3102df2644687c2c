"""Golden fixtures (written by tests/golden/make_golden.py from the reference solver)."""
from __future__ import annotations

import os
from typing import Dict, Tuple

import numpy as np

from golden.make_golden import decks

HERE = os.path.dirname(os.path.abspath(__file__))


def load_golden(name: str):
    deck, steps, keep = decks()[name]
    z = np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False)
    states: Dict[int, Dict[str, np.ndarray]] = {}
    for key in z.files:
        s, field = key.split("/", 1)
        states.setdefault(int(s[1:]), {})[field] = z[key]
    return deck, states


def bitequal(a: np.ndarray, b: np.ndarray) -> bool:
    """Bit-for-bit equality; NaNs compare equal to NaNs (their payloads are platform specific)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype == np.float64:
        na, nb = np.isnan(a), np.isnan(b)
        if not np.array_equal(na, nb):
            return False
        return bool(np.array_equal(np.where(na, 0.0, a).view(np.int64), np.where(nb, 0.0, b).view(np.int64)))
    return bool(np.array_equal(a, b))


def flush_denormal(a: np.ndarray) -> np.ndarray:
    """The reference's FOCC plastic strain rate carries run-to-run-varying denormal residues
    (|x| < 1e-300) from uninitialised memory; they are compared as zero."""
    return np.where(np.abs(a) < 1e-300, 0.0, a)

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
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype == np.float64:
        return bool(np.array_equal(a.view(np.int64), b.view(np.int64)))
    return bool(np.array_equal(a, b))

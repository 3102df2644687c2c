"""sphg_kl_synthesize_batch_noalloc (the K-L batch synthesis with every device buffer in a static
array, for callers that cannot allocate) against sphg_kl_synthesize_batch, through the C-ABI:
bit-equal values on two calls in a row, the second smaller (fewer particles and realisations, so
stale contents of the shared array would show), and the capacity error past 256 realisations."""
import ctypes as C

import numpy as np
import pytest

from sphgpy import slope_2d
from sphgpy.capi import KLField, kl_field, kl_synthesize_batch, product_lib

pytestmark = pytest.mark.gpu

SPHG_E_CAPACITY = 103


def _noalloc(f, xp, ty, xi, fill):
    L = product_lib().lib
    fn = L.sphg_kl_synthesize_batch_noalloc
    fn.argtypes = [C.POINTER(KLField), C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int,
                   C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_double)]
    fn.restype = C.c_int
    xp = np.ascontiguousarray(xp, dtype=np.float64)
    ty = np.ascontiguousarray(ty, np.int32)
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    out = np.full((xi.shape[0], xp.shape[0]), np.nan)
    rc = fn(C.byref(f), xp.shape[0], xp.ctypes.data_as(C.POINTER(C.c_double)), ty.ctypes.data_as(C.POINTER(C.c_int)),
            xi.shape[0], xi.ctypes.data_as(C.POINTER(C.c_double)), fill, out.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, out


def test_noalloc_equals_allocating_entry_point_twice():
    rng = np.random.default_rng(5)
    cases = [(slope_2d(8.0, dx=0.1, seed=3), 200), (slope_2d(3.0, dx=0.1, seed=4), 37)]
    assert cases[0][0].n > 2 * 8192 and cases[1][0].n < cases[0][0].n  # several chunks, then fewer
    for deck, nreal in cases:
        f = kl_field(deck, dist=2, mean=10000.0, sd=3000.0, rl=(10.0, 2.0))
        xi = rng.standard_normal((nreal, 200))
        want, _ = kl_synthesize_batch(f, deck.xp, deck.type, xi, fill=-1.0)
        rc, got = _noalloc(f, deck.xp, deck.type, xi, -1.0)
        assert rc == 0
        print(deck.n, nreal, "max |diff|", float(np.abs(got - want).max()))
        assert np.array_equal(got, want)
        assert np.isfinite(got).all()


def test_noalloc_rejects_more_than_256_realisations():
    deck = slope_2d(3.0, dx=0.2, seed=1)
    f = kl_field(deck, dist=1, mean=1.0, sd=0.1, rl=(10.0, 2.0))
    rc, _ = _noalloc(f, deck.xp, deck.type, np.zeros((257, 200)), 0.0)
    assert rc == SPHG_E_CAPACITY

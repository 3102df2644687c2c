"""Neighbour batch widths, the search's list staging depth and the autotuner
(sphg_set_batch_widths, sphg_set_search_stage, sphg_autotune).

Every width set and both staging depths give the same fields and neighbour lists bit for bit (the
summation order is the list order whatever the width, and the depth only decides which entries
pass through LDS before they are stored); the autotuner leaves the state as it found it and picks
choices from the built set.
"""
import numpy as np
import pytest

from fixtures import load_golden, bitequal
from sphgpy.capi import GpuSim

pytestmark = pytest.mark.gpu

FIELDS = ("xp", "vx", "sig", "rho", "eps", "pre", "ax", "strep", "adps", "vsig")
WIDTHS = ([1, 2, 1, 1], [2, 4, 2, 2], [4, 6, 4, 4], [1, 6, 4, 2])


def make(deck, state, precision=0):
    g = GpuSim(deck, precision=precision)
    g.load_state(state)
    g.upload_done()
    g.set_step(1)
    return g


@pytest.mark.parametrize("name", ["soilwater3d_small", "model_zoo_2d", "tran_2d_b14", "ssi_3d_t1"])
def test_every_width_gives_the_same_results(name):
    deck, states = load_golden(name)
    ref = make(deck, states[0])
    assert ref.step(4) == 0
    for w, stage in zip(WIDTHS, (16, 32, 16, 32)):
        g = make(deck, states[0])
        assert g.set_batch_widths(w) == 0
        assert g.set_search_stage(stage) == 0 and g.search_stage() == stage
        assert g.step(4) == 0
        for k in FIELDS:
            assert bitequal(g.get(k), ref.get(k)), (w, k)
        ca, ia = g.neighbors()
        cb, ib = ref.neighbors()
        assert np.array_equal(ca, cb) and np.array_equal(ia, ib)


def test_autotune_restores_state_and_picks_built_widths():
    deck, states = load_golden("soilwater3d_small")
    g = make(deck, states[0])
    assert g.step(2) == 0
    before = {k: g.get(k) for k in FIELDS}
    step = g.L.lib.sphg_current_step(g.ctx)
    rc, w = g.autotune(3)
    assert rc == 0
    assert w[0] in (1, 2, 4) and w[1] in (2, 4, 6) and w[2] in (1, 2, 4) and w[3] in (1, 2, 4)
    assert g.L.lib.sphg_current_step(g.ctx) == step
    for k in FIELDS:
        assert bitequal(g.get(k), before[k]), k
    ref = make(deck, states[0])
    assert ref.step(2) == 0
    assert g.step(3) == 0 and ref.step(3) == 0
    for k in FIELDS:
        assert bitequal(g.get(k), ref.get(k)), k
    assert g.set_batch_widths([3, 2, 2, 2]) == 100
    assert g.search_stage() in (16, 32)
    assert g.set_search_stage(24) == 100

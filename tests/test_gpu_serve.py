"""Serving loop (sphgpy/serve.py): requests of different sizes admitted while others run, their
work overlapping on the device; every result equals the same request run alone, bit for bit
(each context is deterministic and contexts share no state)."""
import numpy as np
import pytest

from fixtures import bitequal
from sphgpy.serve import Server, request_stream, run_alone

pytestmark = pytest.mark.gpu


def test_results_same_alone_and_in_the_mix():
    reqs = request_stream(6, rate=200.0, seed=3, steps=(20, 60), scale=0.05)
    alone = {r.rid: run_alone(r) for r in reqs}
    srv = Server(hbm_budget=8e9, chunk=7, max_active=6)
    stats = srv.run(reqs)
    assert stats["requests"] == 6 and stats["failed"] == 0
    # requests overlapped: some were admitted before an earlier one finished
    order = sorted(reqs, key=lambda r: r.admitted)
    assert any(b.admitted < a.finished for a, b in zip(order, order[1:]))
    for r in reqs:
        for k, v in alone[r.rid].items():
            assert bitequal(r.result[k], v), (r.rid, r.kind, k)


def test_admission_waits_for_memory():
    reqs = request_stream(4, rate=1000.0, seed=4, steps=(10, 20), scale=0.05)
    srv = Server(hbm_budget=1.0, chunk=5, max_active=4)  # room for one request at a time
    stats = srv.run(reqs)
    assert stats["failed"] == 0
    order = sorted(reqs, key=lambda r: r.admitted)
    assert all(b.admitted >= a.finished for a, b in zip(order, order[1:]))


def test_real_sizes_all_four_families_same_alone_and_in_the_mix():
    # one request of each family at its real size (2D column 5k-50k, 2D random-field slope
    # 20k-80k, 3D soil-water column 100k-500k, 3D slope with pile 0.5M-2M), few steps each,
    # arriving together so that all four run at once on the device
    reqs = request_stream(4, rate=1000.0, seed=7, steps=(3, 8), cycle=True)
    assert sorted(r.kind for r in reqs) == ["column2d", "slope2d", "slopepile3d", "soilwater3d"]
    alone = {r.rid: run_alone(r) for r in reqs}
    stats = Server(hbm_budget=64e9, chunk=2, max_active=4).run(reqs)
    assert stats["requests"] == 4 and stats["failed"] == 0
    print({r.kind: r.n for r in reqs}, stats)
    order = sorted(reqs, key=lambda r: r.admitted)
    assert any(b.admitted < a.finished for a, b in zip(order, order[1:]))
    for r in reqs:
        for k, v in alone[r.rid].items():
            assert bitequal(r.result[k], v), (r.rid, r.kind, k)

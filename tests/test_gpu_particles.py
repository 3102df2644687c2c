"""sphg_upload_particles / sphg_download_particles: the reference's Particle records
(Class_particle.h:13-103, mirrored by sphg_particle, 1064 bytes) to and from the device, against the
per-member sphg_set_field / sphg_get_field path: the same device state (every field bit-equal after
three steps), and downloads equal to the per-member reads, members the device does not keep left
as they were."""
import ctypes as C

import numpy as np
import pytest

from fixtures import load_golden, bitequal
from sphgpy.capi import GpuSim, INT_FIELDS

pytestmark = pytest.mark.gpu

PARTICLE = np.dtype([
    ("x", "<f8", 3), ("xp", "<f8", 3), ("vx", "<f8", 3), ("vxp", "<f8", 3), ("ax", "<f8", 3), ("axp", "<f8", 3),
    ("ux", "<f8", 3), ("strep", "<f8", 3), ("divx", "<f8", 9), ("interf", "<f8", 3), ("interfss", "<f8", 3),
    ("rho", "<f8"), ("rhop", "<f8"), ("mass", "<f8"), ("hl", "<f8"), ("pre", "<f8"), ("prep", "<f8"), ("e", "<f8"),
    ("eps", "<f8", 6), ("deps", "<f8", 6), ("sig", "<f8", 6), ("dsig", "<f8", 6), ("ff_sig", "<f8", 3),
    ("weps", "<f8", 9), ("adps", "<f8", 6), ("tssig", "<f8", 6), ("tsqc", "<f8"), ("satu", "<f8"),
    ("dsatu", "<f8"), ("suct", "<f8"), ("cop", "<f8"), ("vsig", "<f8"), ("veps", "<f8"), ("divq", "<f8"),
    ("divr", "<f8"), ("meps", "<f8"), ("roue", "<f8"), ("rous", "<f8"), ("zeta", "<f8"), ("rsta", "<f8"),
    ("prew", "<f8"), ("prea", "<f8"), ("porosity", "<f8"), ("cd", "<f8"), ("poro_cop", "<f8"),
    ("beta", "<f8", 9), ("beta_ts", "<f8", 9), ("type", "<i4"), ("matype", "<i4"), ("permtype", "<i4"),
    ("etype", "<i4")])
SPHG_PART_ALL = 31
NOT_KEPT = ("divx", "dsig", "ff_sig", "rous", "zeta", "rsta", "tssig", "tsqc")


def _lib():
    from sphgpy.capi import product_lib
    L = product_lib().lib
    L.sphg_upload_particles.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    L.sphg_download_particles.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_uint]
    return L


def _sim(deck):
    g = GpuSim(deck)
    return g


@pytest.mark.parametrize("name", ["air_water_2d_a0", "soilwater3d_small"])
def test_particle_records_match_the_per_field_path(name):
    assert PARTICLE.itemsize == 1064
    L = _lib()
    deck, states = load_golden(name)
    st = states[0]
    a = _sim(deck)
    a.load_state(st)
    a.upload_done()
    a.set_step(1)
    rec = np.zeros(deck.n, PARTICLE)
    for k in PARTICLE.names:
        if k in st:
            rec[k] = np.asarray(st[k]).reshape(rec[k].shape)
    b = _sim(deck)
    assert L.sphg_upload_particles(b.ctx, rec.ctypes.data, deck.n) == 0
    for k, v in st.items():
        if k in ("ps_fai", "ps_c", "ps_cop", "ps_ds"):
            b.set(k, v)
    b.upload_done()
    b.set_step(1)
    assert a.step(3) == 0 and b.step(3) == 0
    kept = [k for k in PARTICLE.names if k not in NOT_KEPT and k not in INT_FIELDS]
    for k in kept:
        assert bitequal(a.get(k), b.get(k)), k
    out = np.zeros(deck.n, PARTICLE)
    for k in NOT_KEPT:
        out[k] = 7.0
    assert L.sphg_download_particles(b.ctx, out.ctypes.data, deck.n, SPHG_PART_ALL) == 0
    for k in kept:
        want = b.get(k)
        assert np.array_equal(out[k].reshape(want.shape).view(np.int64), want.view(np.int64)), k
    for k in INT_FIELDS:
        assert np.array_equal(out[k], b.get(k))
    for k in NOT_KEPT:
        assert np.all(out[k] == 7.0), k
    # a mask downloads only its group
    part = np.zeros(deck.n, PARTICLE)
    assert L.sphg_download_particles(b.ctx, part.ctypes.data, deck.n, 1) == 0  # kinematics
    assert np.array_equal(part["vx"], b.get("vx")) and np.all(part["sig"] == 0.0)
    assert L.sphg_download_particles(b.ctx, part.ctypes.data, deck.n, 64) != 0  # unknown group

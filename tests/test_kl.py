"""K-L random fields (CPP_17-1_K-L_expansion_non.cpp): oracle pinned to the reference, the
library's host-side modes pinned to the oracle, closed forms of the expansion, and (GPU) the
device synthesis against the oracle.

Bars: roots, eigenvalues, mode order and FP64 field values bit-exact between the reference, the
oracle and the library's host code (same C library); device synthesis (OCML cos/sin/log/exp)
within 1e-10 field-max relative; the FP32 matrix-core batch within 2e-5.
"""
import os

import numpy as np
import pytest

from fixtures import bitequal
from helpers import GOLDEN, have_ref, oracle_kl, run_ref_kl, max_rel
from golden.make_golden import kl_cases, kl_inputs

KL = ["kl_slope2d", "kl_block3d"]


def _golden(name):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", KL)
def test_kl_oracle_matches_golden(name):
    g = _golden(name)
    ndim, dist, field = (int(v) for v in g["ints"])
    mean, sd, cdep, smax, wmax = (float(v) for v in g["scal"])
    wi, ev, od, val, rc = oracle_kl(ndim, g["xp"], g["type"], g["a"], g["rl"], g["x_trans"], g["xi"], dist, mean, sd,
                                    cdep, smax, wmax, field, g["value0"])
    assert rc == 0
    assert bitequal(wi, g["wi"]) and bitequal(ev, g["eigval"]) and np.array_equal(od, g["order"])
    assert bitequal(val, g["value"])
    soil = g["type"] == (2 if field <= 3 else 1)
    assert np.std(val[soil]) > 0.05 * abs(mean)  # the field varies


@pytest.mark.skipif(not have_ref(), reason="reference link harness not built")
@pytest.mark.parametrize("name", KL)
def test_kl_oracle_matches_reference(name):
    _, _, args = kl_inputs(name)
    r = run_ref_kl(*args)
    o = oracle_kl(*args)
    assert r[4] == 0 and o[4] == 0
    for a, b in zip(r[:4], o[:4]):
        assert bitequal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("name", KL)
def test_kl_library_modes_match_oracle(name):
    # sphg_kl_domain / sphg_kl_modes are host code: no device needed
    deck, f, args = kl_inputs(name)
    wi, ev, od = f.modes()
    g = _golden(name)
    assert bitequal(np.array(list(f.a)), g["a"]) and bitequal(np.array(list(f.x_trans)), g["x_trans"])
    assert bitequal(wi, g["wi"]) and bitequal(ev, g["eigval"]) and np.array_equal(od, g["order"])


def test_kl_closed_forms():
    g = _golden("kl_slope2d")
    a, rl, wi = g["a"], g["rl"], g["wi"]
    for d in range(2):
        w, aa, dl = wi[:, d], a[d], rl[d]
        i = np.arange(200)
        even = (i + 1) % 2 == 1  # 0-based even modes: 2/delta - w tan(w a) = 0
        # each w is the root of its mode's equation in its own branch of tan (brentq to full
        # precision on the same bracket); the reference stops one halving after |f| <= 1e-6
        from scipy.optimize import brentq
        pi = 3.141592653
        for k in range(200):
            if even[k]:
                lo, hi = k * pi / aa + 1e-6, (k + 0.5) * pi / aa - 1e-6
                root = brentq(lambda x: 2.0 / dl - x * np.tan(x * aa), lo, hi, xtol=1e-15)
            else:
                lo, hi = (k + 0.5) * pi / aa + 1e-6, (k + 1.0) * pi / aa - 1e-6
                root = brentq(lambda x: 2.0 / dl * np.tan(x * aa) + x, lo, hi, xtol=1e-15)
            assert abs(w[k] - root) <= 1e-6 * max(1.0, root), (d, k)
        # eigenfunctions orthonormal on [-a, a] (Gauss-Legendre quadrature)
        xq, wq = np.polynomial.legendre.leggauss(4000)
        x = xq * aa
        alpha = np.where(even, 1.0 / np.sqrt(np.sin(2 * w * aa) / 2 / w + aa),
                         1.0 / np.sqrt(-np.sin(2 * w * aa) / 2 / w + aa))
        phi = np.where(even[:, None], np.cos(np.outer(w, x)), np.sin(np.outer(w, x))) * alpha[:, None]
        gram = (phi * wq * aa) @ phi[:40].T
        assert np.max(np.abs(gram[:40] - np.eye(40))) < 1e-5  # roots carry the 1e-6 bisection tolerance
        # eigenvalues of exp(-2|t|/delta) on [-a, a] sum to the trace 2a over all modes. The reference's
        # brackets take the cosine roots of every other branch of tan (k = 0, 2, 4, ...) and the sine
        # roots of the branches between them (solve_wi, CPP_17-1_K-L_expansion_non.cpp:285-296), i.e.
        # about half of the spectrum: the 200 modes carry just under half of the trace
        lam = 4.0 * dl / (4.0 + w * w * dl * dl)
        assert 0.45 * 2 * aa < lam.sum() < 0.5 * 2 * aa + lam[0]
        assert np.all(np.diff(w) > 0)
    # modes taken in descending eigenvalue order
    assert np.all(np.diff(g["eigval"]) <= 0)


def test_kl_bisection_that_cannot_converge_is_an_error_not_a_hang():
    from sphgpy.capi import KLField, product_lib
    import ctypes as C
    f = KLField()
    f.ndim, f.dist, f.field = 3, 1, 0
    for k, (aa, r) in enumerate(((0.012, 0.02), (0.012, 0.03), (0.0125, 0.01))):
        f.a[k], f.rl[k] = aa, r
    rc = product_lib().lib.sphg_kl_modes(C.byref(f))
    assert rc == 104


# ---------------------------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_kl_device_synthesis_vs_oracle():
    from sphgpy.capi import GpuSim, setting_par
    deck, f, args = kl_inputs("kl_slope2d")
    st = setting_par(deck)
    gsim = GpuSim(deck)
    gsim.load_state(st)
    gsim.upload_done()
    gsim.kl_synthesize(f, args[6])
    got = gsim.get("ps_c")
    *_, want, rc = oracle_kl(*args)
    assert rc == 0
    soil = deck.type == 2
    assert max_rel(got[soil], want[soil]) <= 1e-10
    assert np.std(got[soil]) > 500.0


@pytest.mark.gpu
def test_kl_batch_matrix_core_vs_oracle():
    from sphgpy.capi import kl_synthesize_batch
    deck, f, args = kl_inputs("kl_slope2d")
    rng = np.random.default_rng(7)
    nreal = 40  # not a multiple of the 32-wide tile: exercises the padded columns
    xi = rng.standard_normal((nreal, 200))
    vals, ms = kl_synthesize_batch(f, deck.xp, deck.type, xi, fill=-1.0)
    soil = deck.type == 2
    for r in (0, 17, nreal - 1):
        a = list(args)
        a[6] = xi[r]
        *_, want, rc = oracle_kl(*a)
        assert max_rel(vals[r][soil], want[soil]) <= 2e-5, r
        assert np.all(vals[r][~soil] == -1.0)
    assert ms > 0.0


@pytest.mark.gpu
def test_dp_steps_with_kl_cohesion_bit_exact():
    # the synthesized cohesion enters the Drucker-Prager yield function; with the same per-particle
    # cohesion the step stays bit-exact against the oracle
    from sphgpy.capi import GpuSim, setting_par
    from helpers import OracleSim, JACOBI
    deck, f, args = kl_inputs("kl_slope2d")
    st = setting_par(deck)
    gsim = GpuSim(deck)
    gsim.load_state(st)
    gsim.upload_done()
    gsim.kl_synthesize(f, args[6])
    gsim.set_step(1)
    st2 = dict(st)
    st2["ps_c"] = gsim.get("ps_c")
    o = OracleSim(deck, st2, mode=JACOBI)
    assert gsim.step(5) == 0 and o.step(5) == 0
    for k in ("xp", "vx", "sig", "rho", "eps"):
        assert bitequal(gsim.get(k), o.get(k)), k

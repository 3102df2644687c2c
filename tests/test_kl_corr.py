"""K-L generators and homogeneous random variables of the reference's flspv 1, 2 and 5
(CPP_17-1_K-L_expansion_non.cpp:63-265, CPP_17-2_K-L_expansion_corr.cpp:55-473,
CPP_17-5-Nonstationary-Correlated-RF.cpp:738-863): the oracle against the reference's whole
generators (fixtures tests/golden/kl_spv.npz, written through the link harness), the parts its
clock-seeded normals hide pinned by closed forms, and the library's host code against the oracle.

Bars: bit-exact against the reference where its normals drop out (sd 0; the lognormal branch's 0 / 0
coefficient leaves NaN in the properties fed by the mixed normal, pattern included); the mixing and
the value maps to 1e-15 of their closed forms; the library's host functions bit-exact against the
oracle.
"""
import os

import numpy as np
import pytest

from fixtures import bitequal
from helpers import GOLDEN, have_ref, oracle_kl, oracle_kl_corr_normals, oracle_random_variable, run_ref_rf
from golden.make_rf import KLC_CASES, KLN_CASES, RV_CASES, spv_decks, spv_inputs

RL = (1.0, 0.5, 0.7)


def _fix(name):
    z = np.load(os.path.join(GOLDEN, "kl_spv.npz"), allow_pickle=False)
    return {k.split("/", 1)[1]: z[k] for k in z.files if k.startswith(name + "/")}


def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def kl_domain(d, box):
    """The K-L domain of the reference's generators (CPP_17-1_K-L_expansion_non.cpp:72-90,
    CPP_17-2_K-L_expansion_corr.cpp:71-82): the non-boundary box widened by 20 % per side."""
    a, xt = np.zeros(3), np.zeros(3)
    for k in range(d.ndim):
        lo, hi = box[2 * k], box[2 * k + 1]
        xmin, xmax = lo - 0.2 * (hi - lo), hi + 0.2 * (hi - lo)
        xt[k] = (xmax + xmin) / 2.0
        a[k] = (xmax - xmin) / 2.0
    return a, xt


def _oracle_field(d, box, smax3, wmax3, field, dist, mean, sd, cdep, xi, value):
    a, xt = kl_domain(d, box)
    v = 1 if d.ndim == 2 else 2
    *_, val, rc = oracle_kl(d.ndim, d.xp, d.type, a, RL, xt, xi, dist, mean, sd, cdep, smax3[v], wmax3[v], field,
                            value)
    assert rc == 0
    return val


# ------------------------------------------------------------------ against the reference (sd 0)
@pytest.mark.parametrize("name", sorted(KLN_CASES))
def test_oracle_kl_generator_is_the_references_at_sd_zero(name):
    # flspv 1: every flagged property from its own normals; at sd 0 the normal branch is the depth
    # trend and the lognormal branch exp(log(trend)), the rest untouched
    z = _fix(name)
    dk, spec = KLN_CASES[name]
    d = spv_decks()[dk]
    box, smax3, wmax3, init = spv_inputs(d)
    assert int(z["rc"][0]) == 0
    got = np.array(init)
    xi = np.random.default_rng(1).standard_normal(200)
    for fld, (dist, mean, cdep) in spec.items():
        got[fld] = _oracle_field(d, box, smax3, wmax3, fld, dist, mean, 0.0, cdep, xi, got[fld])
    assert np.array_equal(_bits(got), _bits(z["value"]))
    for fld in range(6):
        assert np.array_equal(got[fld], z["init"][fld]) == (fld not in spec)


@pytest.mark.parametrize("name", sorted(KLC_CASES))
def test_oracle_kl_corr_generator_is_the_references_at_sd_zero(name):
    # flspv 2: the pair (fai, c) of each flagged phase from mixed normals, each with its own
    # distribution; a lognormal fai makes the coefficient 0 / 0 at sd 0 and the c field NaN
    z = _fix(name)
    dk, spec, coe = KLC_CASES[name]
    d = spv_decks()[dk]
    box, smax3, wmax3, init = spv_inputs(d)
    assert int(z["rc"][0]) == 0
    got = np.array(init)
    u = np.random.default_rng(2).standard_normal((200, 2))
    for fld, (pa, pb) in spec.items():
        xa, xb, _ = oracle_kl_corr_normals(pa[0], pa[1], 0.0, pb[1], 0.0, coe, u)
        got[fld] = _oracle_field(d, box, smax3, wmax3, fld, pa[0], pa[1], 0.0, pa[2], xa, got[fld])
        got[fld + 1] = _oracle_field(d, box, smax3, wmax3, fld + 1, pb[0], pb[1], 0.0, pb[2], xb, got[fld + 1])
        assert np.isnan(got[fld + 1]).any() == (pa[0] == 2)
    assert np.array_equal(_bits(got), _bits(z["value"]))


@pytest.mark.parametrize("name", sorted(KLN_CASES) + sorted(KLC_CASES))
def test_kl_domain_is_the_references(name):
    # the half-widths a the reference's generators set for every property, against the library's
    # sphg_kl_domain (host code) and this file's restatement
    from sphgpy.capi import kl_field
    z = _fix(name)
    dk = (KLN_CASES.get(name) or KLC_CASES[name])[0]
    d = spv_decks()[dk]
    box, *_ = spv_inputs(d)
    a, _ = kl_domain(d, box)
    f = kl_field(d, 1, 10.0, 1.0, RL)
    for k in range(6):
        assert bitequal(z["a"][k][:d.ndim], a[:d.ndim])
        assert bitequal(np.array(list(f.a))[:d.ndim], a[:d.ndim])
    _, xt = kl_domain(d, box)
    assert bitequal(np.array(list(f.x_trans))[:d.ndim], xt[:d.ndim])


@pytest.mark.parametrize("name", sorted(RV_CASES))
def test_oracle_random_variable_is_the_references_at_sd_zero(name):
    # flspv 5: one value per flagged property for the whole phase; the distribution comes from
    # psv[0] (flag_fai), else psv[4] (flag_faiw), else the constructor's normal
    z = _fix(name)
    dk, flags, spec, coe = RV_CASES[name]
    d = spv_decks()[dk]
    _, _, _, init = spv_inputs(d)
    assert int(z["rc"][0]) == 0
    psv = [(1, 0.0, 0.0)] * 6
    for fld, (dist, mean) in spec.items():
        psv[fld] = (dist, mean, 0.0)
    got = oracle_random_variable(flags, coe, psv, (0.37, -1.21), d.type, init)
    assert np.array_equal(_bits(got), _bits(z["value"]))


def test_fixture_cases_cover_the_branches():
    z = {n: _fix(n) for n in list(KLN_CASES) + list(KLC_CASES) + list(RV_CASES)}
    # lognormal pairs leave NaN exactly in the property of the mixed normal
    assert np.isnan(z["klc_box3d_soil_lognormal"]["value"][1]).any()
    assert not np.isnan(z["klc_box3d_soil_lognormal"]["value"][0]).any()
    d = spv_decks()["box2d"]
    soil = d.type == 2
    assert np.isnan(z["rv_soil_lognormal"]["value"][0][soil]).all()
    assert np.isfinite(z["rv_soil_lognormal"]["value"][1]).all()
    # flag_c alone reads the default (normal) distribution, not psv[1]'s lognormal one
    assert np.all(z["rv_c_only_default_normal"]["value"][1][soil] == 9000.3)
    assert np.exp(np.log(9000.3)) != 9000.3  # so a lognormal map would have shown


# ------------------------------------------------------------------ closed forms
def _lognormal_corr(rho_n, cv_a, cv_b):
    """Correlation of exp(A), exp(B) for normals A, B with correlation rho_n and the variances of
    lognormals with coefficients of variation cv_a, cv_b."""
    sa, sb = np.sqrt(np.log1p(cv_a ** 2)), np.sqrt(np.log1p(cv_b ** 2))
    return np.expm1(rho_n * sa * sb) / np.sqrt(np.expm1(sa * sa) * np.expm1(sb * sb))


@pytest.mark.parametrize("coe", [-0.7, -0.2, 0.3, 0.8])
def test_lognormal_coefficient_gives_the_asked_correlation(coe):
    # the coefficient of the underlying normals, mapped forward, is the asked correlation of the
    # lognormal properties
    u = np.zeros((200, 2))
    _, _, rho = oracle_kl_corr_normals(2, 25.0, 6.0, 9000.0, 2700.0, coe, u)
    assert abs(_lognormal_corr(rho, 6.0 / 25.0, 2700.0 / 9000.0) - coe) <= 1e-12
    _, _, rho1 = oracle_kl_corr_normals(1, 25.0, 6.0, 9000.0, 2700.0, coe, u)
    assert rho1 == coe


def test_mixing_limits_and_linearity():
    u = np.random.default_rng(3).standard_normal((200, 2))
    xa, xb, _ = oracle_kl_corr_normals(1, 1.0, 1.0, 1.0, 1.0, 0.0, u)
    assert np.array_equal(xa, u[:, 0]) and np.array_equal(xb, u[:, 1])
    xa, xb, _ = oracle_kl_corr_normals(1, 1.0, 1.0, 1.0, 1.0, 1.0, u)
    assert np.array_equal(xa, u[:, 0]) and np.array_equal(xb, u[:, 0])
    xa, xb, _ = oracle_kl_corr_normals(1, 1.0, 1.0, 1.0, 1.0, 0.6, u)
    assert np.array_equal(xa, u[:, 0])
    assert np.allclose(xb, 0.6 * u[:, 0] + 0.8 * u[:, 1], rtol=0, atol=1e-15)


def test_random_variable_value_maps():
    ty = np.array([2, 1, 2, 0, 3, 1], np.int32)
    init = np.full((6, 6), -1.0)
    z = (0.9, -0.4)
    coe = 0.6
    # normal: c takes the first normal, fai the mixed one
    v = oracle_random_variable((1, 1, 0, 0), coe, [(1, 30.0, 3.0), (1, 800.0, 50.0)] + [(1, 0.0, 0.0)] * 4, z, ty,
                               init)
    assert v[0][0] == pytest.approx(30.0 + 3.0 * (0.6 * 0.9 + 0.8 * -0.4), rel=1e-15)
    assert v[1][0] == 800.0 + 50.0 * 0.9
    assert np.array_equal(v[0][ty != 2], init[0][ty != 2]) and np.all(v[0][ty == 2] == v[0][0])
    # lognormal: at zero normals the value is the median mean / sqrt(1 + cv^2)
    v = oracle_random_variable((0, 0, 1, 1), coe, [(1, 0.0, 0.0)] * 4 + [(2, 4.0, 1.0), (2, 2.0, 0.4)], (0.0, 0.0),
                               ty, init)
    assert v[4][1] == pytest.approx(4.0 / np.sqrt(1.0 + 0.25 ** 2), rel=1e-14)
    assert v[5][1] == pytest.approx(2.0 / np.sqrt(1.0 + 0.2 ** 2), rel=1e-14)
    assert np.all(v[4][ty == 3] == -1.0)  # air keeps its value: only water (type 1) is written


# ------------------------------------------------------------------ library host code
def test_library_corr_normals_equal_the_oracle():
    from sphgpy.capi import KLField, kl_corr_normals
    u = np.random.default_rng(5).standard_normal((200, 2))
    for dist, coe in ((1, 0.35), (2, -0.45), (2, 0.0)):
        fa, fb = KLField(), KLField()
        fa.dist, fa.mean_value, fa.sd_value = dist, 25.0, 5.0
        fb.dist, fb.mean_value, fb.sd_value = 1, 9000.0, 2000.0
        xa, xb, rho = kl_corr_normals(fa, fb, coe, u)
        oa, ob, orho = oracle_kl_corr_normals(dist, 25.0, 5.0, 9000.0, 2000.0, coe, u)
        assert bitequal(xa, oa) and bitequal(xb, ob) and rho == orho


@pytest.mark.parametrize("flags", [(1, 1, 0, 0), (0, 0, 1, 1), (1, 1, 1, 1), (0, 1, 0, 0), (0, 0, 1, 0)])
@pytest.mark.parametrize("dists", [(1, 1), (2, 2), (2, 1)])
def test_library_random_variable_values_equal_the_oracle(flags, dists):
    from sphgpy.capi import random_variable_values
    psv = [(dists[0], 31.0, 4.0), (dists[0], 850.0, 120.0), (1, 0.0, 0.0), (1, 0.0, 0.0), (dists[1], 6.0, 1.5),
           (dists[1], 1.7, 0.3)]
    z = (-0.8, 1.3)
    v = random_variable_values(flags, 0.4, psv, z)
    ty = np.array([2, 1], np.int32)
    o = oracle_random_variable(flags, 0.4, psv, z, ty, np.full((6, 2), np.nan))
    want = np.array([o[0][0], o[1][0], np.nan, np.nan, o[4][1], o[5][1]])
    assert np.array_equal(_bits(v), _bits(want))


@pytest.mark.skipif(not have_ref(), reason="reference link harness not built")
def test_reference_random_variable_live():
    # the live reference at sd 0 on another deck and coefficient than the fixtures
    d = spv_decks()["slope"]
    box, smax3, wmax3, init = spv_inputs(d)
    psv = [(1, 0.0, 0.0, RL, 0.0)] * 6
    psv[0] = (2, 27.0, 0.0, RL, 0.0)
    psv[1] = (2, 12345.6, 0.0, RL, 0.0)
    vals, rc = run_ref_rf(d.ndim, d.xp, d.type, d.dr, box, smax3, wmax3, [1, 1, 0, 0, 0, 0], 1.0, 1, 0.0, 0.0, psv,
                          init, coe=-0.35, kind="rv")
    assert rc == 0
    got = oracle_random_variable((1, 1, 0, 0), -0.35, [p[:3] for p in psv], (0.1, 0.2), d.type, init)
    assert np.array_equal(_bits(got), _bits(vals))


# ------------------------------------------------------------------ Parameters.txt (flspv -1)
def _parameters_text(d, vals, with_boundary_lines):
    """A Parameters.txt in the layout of the reference's writer (CPP_18-spatialoutput.cpp:34-52),
    which lists water, soil, air and structure particles only; with_boundary_lines adds a line for
    every other particle too."""
    rows = ["No.---- X----Y----Z----fai----c----cop----ds----type----matype----"]
    n = d.n
    for i in range(n):
        t = int(d.type[i])
        if t in (2, 4):
            v = (vals[0][i], vals[1][i], vals[2][i], vals[3][i])
        elif t in (1, 3):
            v = (vals[4][i], vals[5][i], vals[2][i], vals[3][i])
        elif with_boundary_lines:
            v = (0.0, 0.0, 0.0, 0.0)
        else:
            continue
        x = d.xp[i]
        rows.append("   %6d   %10.6e   %10.6e   %10.6e   %10.6e   %10.6e   %10.6e   %10.6e   %2d   %2d"
                    % (i + 1, x[0], x[1], x[2], *v, t, int(d.matype[i])))
    rows.append(rows[0])
    return "\n".join(rows) + "\n"


@pytest.mark.parametrize("with_boundary_lines", [True, False])
def test_parameters_file_reader(with_boundary_lines, tmp_path):
    # the library's reader against the parse of the written values and, where the reference is
    # built, against its own random_field_input; the box lists its walls last, so a file without
    # boundary lines reads the same for every water and soil particle
    import ctypes as C
    from sphgpy.capi import product_lib
    d = spv_decks()["box3d"]
    rng = np.random.default_rng(12)
    vals = rng.uniform(0.5, 50.0, (6, d.n))
    text = _parameters_text(d, vals, with_boundary_lines)
    (tmp_path / "Parameters.txt").write_text(text)
    init = rng.uniform(-2.0, -1.0, (6, d.n))
    got = np.ascontiguousarray(init.copy())
    ty = np.ascontiguousarray(d.type, np.int32)
    rc = product_lib().lib.sphg_read_parameters(str(tmp_path).encode(), d.n, ty.ctypes.data_as(C.POINTER(C.c_int)),
                                                got.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0
    printed = np.array([[float("%10.6e" % v) for v in row] for row in vals])
    for k, phase in ((0, (2, 4)), (1, (2, 4)), (2, (2, 4)), (3, (2, 4)), (4, (1, 3)), (5, (1, 3))):
        m = np.isin(d.type, phase)
        assert np.array_equal(got[k][m], printed[k][m]), k
        assert np.array_equal(got[k][~m], init[k][~m]), k
    if have_ref():
        _, smax3, wmax3, _ = spv_inputs(d)
        ref, rc_ref = run_ref_rf(d.ndim, d.xp, d.type, d.dr, [0.0] * 6, smax3, wmax3, [0] * 6, 1.0, 1, 0.0, 0.0,
                                 [(1, 0.0, 0.0, RL, 0.0)] * 6, list(init), kind="rfi",
                                 files={"Parameters.txt": text})
        assert rc_ref == 0
        assert np.array_equal(_bits(got), _bits(ref))
    assert product_lib().lib.sphg_read_parameters(str(tmp_path / "none").encode(), d.n,
                                                   ty.ctypes.data_as(C.POINTER(C.c_int)),
                                                   got.ctypes.data_as(C.POINTER(C.c_double))) == 100


@pytest.mark.skipif(not have_ref(), reason="reference link harness not built")
def test_parameters_file_with_leading_walls_reads_like_the_reference(tmp_path):
    # walls first: a file in the writer's layout (no wall lines) is read one line per wall, so the
    # values land on later particles than they were written for - in the reference and here alike
    import ctypes as C
    from sphgpy.capi import product_lib
    d0 = spv_decks()["slope"]
    perm = np.argsort(d0.type != 0, kind="stable")  # walls first, then the soil in its order
    d = d0.copy()
    for k in ("xp", "vxp", "type", "matype", "permtype_p", "etype"):
        setattr(d, k, getattr(d0, k)[perm].copy())
    assert d.type[0] == 0 and d.type[-1] == 2
    rng = np.random.default_rng(13)
    vals = rng.uniform(0.5, 50.0, (6, d.n))
    text = _parameters_text(d, vals, False)
    (tmp_path / "Parameters.txt").write_text(text)
    init = rng.uniform(-2.0, -1.0, (6, d.n))
    got = np.ascontiguousarray(init.copy())
    ty = np.ascontiguousarray(d.type, np.int32)
    assert product_lib().lib.sphg_read_parameters(str(tmp_path).encode(), d.n, ty.ctypes.data_as(C.POINTER(C.c_int)),
                                                  got.ctypes.data_as(C.POINTER(C.c_double))) == 0
    _, smax3, wmax3, _ = spv_inputs(d)
    ref, rc_ref = run_ref_rf(d.ndim, d.xp, d.type, d.dr, [0.0] * 6, smax3, wmax3, [0] * 6, 1.0, 1, 0.0, 0.0,
                             [(1, 0.0, 0.0, RL, 0.0)] * 6, list(init), kind="rfi", files={"Parameters.txt": text})
    assert rc_ref == 0
    assert np.array_equal(_bits(got), _bits(ref))
    soil = d.type == 2
    printed = np.array([float("%10.6e" % v) for v in vals[0]])
    assert not np.array_equal(got[0][soil], printed[soil])  # shifted, as in the reference

"""Correlation-matrix random fields (flspv 3): the oracle against the reference, closed forms and
invariants, and the library's host cell partition against the oracle. Host only.

Pins of the oracle (oracle/rf_oracle.c):
- Cholesky: bit-equal to the reference's own Cholesky (fixtures from `ref_harness --chol`) on
  correlation matrices of five cases, two of them in the clamp regime; and equal to the closed
  form of a Kronecker product of AR(1) factors on a separable exponential field.
- Correlation matrix: symmetric, unit diagonal, the separable forms equal to Kronecker products,
  the rotated form with zero angles equal to the exponential form bit for bit.
- Cell partition and per-particle values: bit-equal to the reference's nonstationary_random_2D/3D
  with sd 0 (where its clock-seeded normals drop out; fixtures from `ref_harness --rf`),
  including its error codes 41 and 42; for sd > 0 the value maps fixed by the mathematics
  (normal: linear in E with slope sd; lognormal: mean of the distribution equal to the trend).
"""
import os

import numpy as np
import pytest

from golden.make_rf import CHOL_CASES, PIPE_CASES, chol_matrix, corr_of, digest, pipe_inputs, rf_decks
from helpers import oracle_rf, oracle_chol
from sphgpy.capi import rf_cells, rf_field

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _chol_fixture():
    return np.load(os.path.join(GOLDEN, "rf_chol.npz"), allow_pickle=False)


def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.int64)


@pytest.mark.parametrize("name", sorted(CHOL_CASES))
def test_oracle_cholesky_is_the_references(name):
    z = _chol_fixture()
    f, d, cells = chol_matrix(name)
    A = corr_of(d, f, cells)
    assert digest(A) == bytes(z[f"{name}/digest"]).decode()  # the matrix the reference factored
    assert np.array_equal(_bits(oracle_chol(A)), _bits(z[f"{name}/L"]))


def test_fixture_cases_cover_the_clamp():
    z = _chol_fixture()
    clamped = {n: int((np.diag(z[f"{n}/L"]) == np.sqrt(1e-6)).sum()) for n in CHOL_CASES}
    assert clamped["exp2d"] == 0 and clamped["rot3d"] == 0
    assert clamped["gauss2d_smooth"] > 0 and clamped["rot2d"] > 0


def ar1_factor(m, rho):
    """Cholesky factor of the AR(1) correlation rho^|i-j|: column 0 is rho^i, column j >= 1 is
    rho^(i-j) sqrt(1 - rho^2)."""
    L = np.zeros((m, m))
    for i in range(m):
        L[i, 0] = rho ** i
        for j in range(1, i + 1):
            L[i, j] = rho ** (i - j) * np.sqrt(1.0 - rho * rho)
    return L


def separable_case(ndim=2):
    d = rf_decks()["slope"]
    f = rf_field(d, 1, 1.0, 0.0, (3.0, 1.2), acf=1, cell_ndx=5.0)
    rc, _, cells, L, _ = oracle_rf(2, d.xp, d.type, f, np.zeros(1), want_l=True)
    assert rc == 0
    return d, f, cells, L


def test_oracle_cholesky_closed_form_on_a_separable_field():
    d, f, cells, L = separable_case()
    nx, ny = int(cells["nc"][0]), int(cells["nc"][1])
    rx = np.exp(-2.0 * cells["cell_dx"] / f.rl[0])
    ry = np.exp(-2.0 * cells["cell_dx"] / f.rl[1])
    closed = np.kron(ar1_factor(ny, ry), ar1_factor(nx, rx))  # cell id j * nx + i
    assert np.abs(L - closed).max() < 1e-12


@pytest.mark.parametrize("acf", [1, 2, 3])
def test_correlation_matrix_invariants(acf):
    d = rf_decks()["slope"]
    f = rf_field(d, 1, 1.0, 0.0, (3.0, 1.2), acf=acf, cell_ndx=5.0, beta=0.0)
    _, _, cells, _, _ = oracle_rf(2, d.xp, d.type, f, np.zeros(1))
    A = corr_of(d, f, cells)
    assert np.array_equal(A, A.T) and np.all(np.diag(A) == 1.0)
    nx, ny = int(cells["nc"][0]), int(cells["nc"][1])
    h = cells["cell_dx"] * np.arange(max(nx, ny))
    if acf in (1, 3):
        cx = np.exp(-2.0 * np.abs(h[:nx, None] - h[None, :nx]) / f.rl[0])
        cy = np.exp(-2.0 * np.abs(h[:ny, None] - h[None, :ny]) / f.rl[1])
    else:
        pi = 3.141592653  # the reference's pi (Header_Parameters.h:10)
        cx = np.exp(-pi * (np.abs(h[:nx, None] - h[None, :nx]) / f.rl[0]) ** 2)
        cy = np.exp(-pi * (np.abs(h[:ny, None] - h[None, :ny]) / f.rl[1]) ** 2)
    assert np.abs(A - np.kron(cy, cx)).max() < 1e-13
    if acf == 3:  # the rotated form at zero angle is the exponential form, bit for bit
        g = rf_field(d, 1, 1.0, 0.0, (3.0, 1.2), acf=1, cell_ndx=5.0)
        assert np.array_equal(A, corr_of(d, g, cells))


def test_rotated_3d_form_at_zero_angles_is_the_exponential_form():
    d = rf_decks()["box3d"]
    f3 = rf_field(d, 1, 1.0, 0.0, (0.008, 0.004, 0.006), acf=3, cell_ndx=1.0)
    f1 = rf_field(d, 1, 1.0, 0.0, (0.008, 0.004, 0.006), acf=1, cell_ndx=1.0)
    _, _, cells, _, _ = oracle_rf(3, d.xp, d.type, f3, np.zeros(1))
    assert np.array_equal(corr_of(d, f3, cells), corr_of(d, f1, cells))


@pytest.mark.parametrize("name", sorted(PIPE_CASES))
def test_oracle_field_is_the_references_at_sd_zero(name):
    z = np.load(os.path.join(GOLDEN, "rf_pipeline.npz"), allow_pickle=False)
    d, f = pipe_inputs(name)
    rc_ref = int(z[f"{name}/rc"][0])
    rng = np.random.default_rng(1)
    rc, value, cells, _, _ = oracle_rf(d.ndim, d.xp, d.type, f, rng.standard_normal(15000),
                                       fill=np.nan)
    assert rc == rc_ref
    # the library's host partition agrees, error codes included
    rcl, cell_of = rf_cells(f, d.xp, d.type)
    assert rcl == rc_ref
    if rc:
        return
    assert np.array_equal(cell_of, cells["cell_of"])
    assert list(f.nc) == [int(v) for v in cells["nc"]] and f.cell_total == cells["total"]
    assert np.array_equal(_bits(list(f.base)), _bits(cells["base"])) and f.cell_dx == cells["cell_dx"]
    got = np.where(np.isnan(value), z[f"{name}/init"], value)
    assert np.array_equal(_bits(got), _bits(z[f"{name}/value"]))
    assert not np.array_equal(got, z[f"{name}/init"])  # the field did change the property


def test_value_maps_with_spread():
    # normal: linear in E with slope sd; lognormal: exp(st^2 / 2) v(E = 0) is the trend mean and
    # log v is linear in E with slope st = sqrt(log(1 + (sd / m)^2))
    d = rf_decks()["slope"]
    for dist in (1, 2):
        f = rf_field(d, dist, 50.0, 10.0, (2.0, 0.6), field=1, cell_ndx=5.0, coe_depth=3.0)
        rc, _, cells, Lm, _ = oracle_rf(2, d.xp, d.type, f, np.zeros(144), want_l=True)
        total = cells["total"]
        zeta = np.linalg.solve(Lm, np.linspace(-2.0, 2.0, total))  # E = L zeta spans [-2, 2]
        rc, v, cells, _, E = oracle_rf(2, d.xp, d.type, f, zeta)
        assert rc == 0 and np.abs(E - np.linspace(-2.0, 2.0, total)).max() < 1e-9
        _, v0, _, _, _ = oracle_rf(2, d.xp, d.type, f, np.zeros(total))
        m = d.type == 2
        c = cells["cell_of"][m]
        e = E[c]
        y = cells["base"][1] + (c // cells["nc"][0] + 0.5) * cells["cell_dx"]
        trend = 50.0 + 3.0 * np.maximum(f.soil_max - y, 0.0) * (f.soil_max > y)
        if dist == 1:
            assert np.abs(v[m] - v0[m] - 10.0 * e).max() < 1e-9 * np.abs(v).max()
            assert np.abs(v0[m] - trend).max() < 1e-12 * trend.max()
        else:
            st = np.sqrt(np.log(1.0 + (10.0 / trend) ** 2))
            assert np.abs(v0[m] * np.exp(st * st / 2) / trend - 1.0).max() < 1e-12
            assert np.abs(np.log(v[m]) - np.log(v0[m]) - st * e).max() < 1e-12


# ------------------------------------------------------------ correlated pair (flspv 4)
from golden.make_rf import PAIR_CASES, pair_inputs  # noqa: E402
from helpers import oracle_rf_pair  # noqa: E402


@pytest.mark.parametrize("name", sorted(PAIR_CASES))
def test_oracle_pair_is_the_references_at_sd_zero(name):
    # nonstationary_correlated_random_2D/3D with sd 0 (normal): the mixing and the normals drop
    # out, the cell partition, the trend axes and tops (the fluid pair on y in 3D too) remain
    z = np.load(os.path.join(GOLDEN, "rf_pair.npz"), allow_pickle=False)
    d, fa, fc, coe = pair_inputs(name)
    assert int(z[f"{name}/rc"][0]) == 0
    rc, va, vc = oracle_rf_pair(d.ndim, d.xp, d.type, fa, fc, coe,
                                np.random.default_rng(2).standard_normal((15000, 2)), fill=np.nan)
    assert rc == 0
    for got, init, want in ((va, "init_fai", "fai"), (vc, "init_c", "c")):
        got = np.where(np.isnan(got), z[f"{name}/{init}"], got)
        assert np.array_equal(_bits(got), _bits(z[f"{name}/{want}"])), want
        assert not np.array_equal(got, z[f"{name}/{init}"])


def test_oracle_pair_mixing():
    # coefficient 1: the cohesion sees the friction angle's normals (equal values for equal
    # parameters); coefficient 0: the cohesion sees the second normals alone
    d, fa, fc, _ = pair_inputs("slope_soil_pair")
    for f in (fa, fc):
        f.sd_value, f.mean_value, f.coe_depth = 3.0, 20.0, 1.0
    rc, _, cells, _, _ = oracle_rf(2, d.xp, d.type, fa, np.zeros(1), cells_only=True)
    z1 = np.random.default_rng(4).standard_normal((cells["total"], 2))
    _, va, vc = oracle_rf_pair(2, d.xp, d.type, fa, fc, 1.0, z1)
    assert np.array_equal(va, vc)
    _, va0, vc0 = oracle_rf_pair(2, d.xp, d.type, fa, fc, 0.0, z1)
    _, v1, _ = oracle_rf_pair(2, d.xp, d.type, fa, fc, 0.0, z1[:, ::-1].copy())
    assert np.array_equal(vc0, v1)

"""Correlation-matrix random fields on the GPU (sphg_rf_generate) against the oracle.

Bars (FP64; the device factor is blocked and sums in another order than the reference's
column-by-column Cholesky, so it is not bit-equal):
- factor of a positive definite correlation matrix (no clamped diagonal): max |L_gpu - L_oracle|
  <= 1e-12 (entries of L are at most 1); closed form of a separable field <= 1e-11 at 1,600 cells;
- where the clamp sqrt(max(d, 1e-6)) acts on a nearly singular matrix the factor is not a
  continuous function of rounding past the first clamped column, so the clamp pattern (which
  diagonals are clamped) must be equal and the columns before the first clamp agree (<= 1e-9);
- an indefinite matrix whose factor degenerates to inf / NaN despite the clamp (the reference then
  assigns NaN properties) is refused with SPHG_E_NONFINITE (105); the oracle's factor of the same
  matrix is confirmed non-finite;
- per-particle values <= 1e-9 field-max relative (device log/pow/exp);
- every realisation of a batch equals that realisation generated alone, bit for bit;
- the reference's error codes 41 and 42.
"""
import numpy as np
import pytest

from golden.make_rf import CHOL_CASES, PIPE_CASES, chol_matrix, pipe_inputs, rf_decks
from helpers import oracle_rf, max_rel
from sphgpy.capi import rf_field, rf_generate

pytestmark = pytest.mark.gpu

CLAMPED = ("gauss2d_smooth",)
NONFINITE = ("rot2d",)
SPHG_E_NONFINITE = 105


@pytest.mark.parametrize("name", sorted(CHOL_CASES))
def test_factor_matches_oracle(name):
    f, d, cells = chol_matrix(name)
    total = cells["total"]
    zeta = np.random.default_rng(3).standard_normal(total)
    rc, value, Lg, ms = rf_generate(f, d.xp, d.type, zeta, want_l=True)
    rco, vo, _, Lo, _ = oracle_rf(d.ndim, d.xp, d.type, f, zeta, want_l=True)
    assert rco == 0
    if name in NONFINITE:
        assert not np.isfinite(Lo).all()
        assert rc == SPHG_E_NONFINITE
        return
    assert rc == 0
    assert np.isfinite(Lg).all() and np.isfinite(Lo).all()
    assert np.all(np.triu(Lg, 1) == 0.0)
    clamp_g = np.diag(Lg) == np.sqrt(1e-6)
    clamp_o = np.diag(Lo) == np.sqrt(1e-6)
    print(name, total, "max |dL|", np.abs(Lg - Lo).max(), "clamped", int(clamp_o.sum()), int(clamp_g.sum()))
    if name in CLAMPED:
        assert clamp_o.any()
        assert np.array_equal(clamp_g, clamp_o)
        first = int(np.argmax(clamp_o))  # before the first clamped column the factors agree
        assert np.abs(Lg[:, :first] - Lo[:, :first]).max() <= 1e-9
        assert np.isfinite(value).all()
        return
    assert not clamp_o.any()
    assert np.abs(Lg - Lo).max() <= 1e-12
    assert max_rel(value[0], vo) <= 1e-9


def test_closed_form_separable_field_at_1600_cells():
    d = rf_decks()["slope"]
    f = rf_field(d, 1, 1.0, 0.0, (3.0, 1.2), acf=1, cell_ndx=1.5)
    from sphgpy.capi import rf_cells
    g = rf_field(d, 1, 1.0, 0.0, (3.0, 1.2), acf=1, cell_ndx=1.5)
    assert rf_cells(g, d.xp, d.type)[0] == 0
    nx, ny = g.nc[0], g.nc[1]
    assert g.cell_total >= 1600
    rc, _, Lg, ms = rf_generate(f, d.xp, d.type, np.zeros(g.cell_total), want_l=True)
    assert rc == 0

    def ar1(m, rho):
        i = np.arange(m)
        L = np.where(i[:, None] >= i[None, :], rho ** np.maximum(i[:, None] - i[None, :], 0), 0.0)
        L[:, 1:] *= np.sqrt(1.0 - rho * rho)
        return L

    closed = np.kron(ar1(ny, np.exp(-2.0 * g.cell_dx / 1.2)), ar1(nx, np.exp(-2.0 * g.cell_dx / 3.0)))
    print("cells", g.cell_total, "max |dL|", np.abs(Lg - closed).max(), "factor ms", ms)
    assert np.abs(Lg - closed).max() <= 1e-11


@pytest.mark.parametrize("name", sorted(PIPE_CASES))
def test_values_match_oracle_with_spread(name):
    d, f = pipe_inputs(name)
    f.sd_value = 0.3 * abs(f.mean_value)
    rng = np.random.default_rng(7)
    zeta = rng.standard_normal((3, 15000))
    rco, vo, cells, _, _ = oracle_rf(d.ndim, d.xp, d.type, f, zeta[1], fill=-7.0)
    total = max(cells["total"], 1)
    rc, v, _, _ = rf_generate(f, d.xp, d.type, np.ascontiguousarray(zeta[:, :total]), fill=-7.0)
    assert rc == rco
    if rc:
        return
    assert max_rel(v[1], vo) <= 1e-9, max_rel(v[1], vo)
    assert np.array_equal(v[1] == -7.0, vo == -7.0)
    # each realisation of the batch is what it is alone
    rc1, v1, _, _ = rf_generate(f, d.xp, d.type, zeta[2, :total], fill=-7.0)
    assert rc1 == 0 and np.array_equal(v1[0], v[2])


# ------------------------------------------------------------ correlated pair (flspv 4)
from golden.make_rf import PAIR_CASES, pair_inputs  # noqa: E402
from helpers import oracle_rf_pair  # noqa: E402
from sphgpy.capi import rf_generate_pair, rf_cells  # noqa: E402


@pytest.mark.parametrize("dist", [1, 2])
@pytest.mark.parametrize("name", sorted(PAIR_CASES))
def test_pair_matches_oracle(name, dist):
    d, fa, fc, coe = pair_inputs(name)
    for f in (fa, fc):
        f.dist = dist
        f.sd_value = 0.25 * f.mean_value
    g = rf_field(d, 1, 1.0, 0.0, (1.0, 0.5, 0.7), field=fa.field, cell_ndx=fa.cell_ndx)
    assert rf_cells(g, d.xp, d.type)[0] == 0
    z1 = np.random.default_rng(11).standard_normal((2, g.cell_total, 2))
    rc, va, vc = rf_generate_pair(fa, fc, coe, d.xp, d.type, z1, fill=-5.0)
    assert rc == 0
    for r in range(2):
        rco, oa, oc = oracle_rf_pair(d.ndim, d.xp, d.type, fa, fc, coe, z1[r], fill=-5.0)
        assert rco == 0
        assert max_rel(va[r], oa) <= 1e-9 and max_rel(vc[r], oc) <= 1e-9, (max_rel(va[r], oa), max_rel(vc[r], oc))


def test_pair_without_correlation_is_two_single_fields():
    # coefficient 0, normal: each property is the single-field generation of its own normals
    d, fa, fc, _ = pair_inputs("slope_soil_pair")
    for f in (fa, fc):
        f.sd_value = 0.3 * f.mean_value
    g = rf_field(d, 1, 1.0, 0.0, (1.0, 0.5), field=0, cell_ndx=fa.cell_ndx)
    assert rf_cells(g, d.xp, d.type)[0] == 0
    z1 = np.random.default_rng(12).standard_normal((g.cell_total, 2))
    rc, va, vc = rf_generate_pair(fa, fc, 0.0, d.xp, d.type, z1)
    assert rc == 0
    rc1, a1, _, _ = rf_generate(fa, d.xp, d.type, np.ascontiguousarray(z1[:, 0]))
    rc2, c1, _, _ = rf_generate(fc, d.xp, d.type, np.ascontiguousarray(z1[:, 1]))
    assert rc1 == 0 and rc2 == 0
    assert np.array_equal(va[0], a1[0]) and np.array_equal(vc[0], c1[0])

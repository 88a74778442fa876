"""Fit statistics summed in the response fold's epilogue: chi2 of the
reference's cal_chi2 (p_devarshi24st.py:43-49, model and data scaled so the
model sums to 1e6) and the Poisson log-likelihood, checked against the same
formulas evaluated with numpy/scipy on the returned counts, for shared and
per-star data, on the batched and the plain path, and against the FP64
oracle's counts."""
import numpy as np
import pytest
from scipy.special import gammaln

from gpu_ppm_amd import abi, fixtures
from gpu_ppm_amd import scenarios as S
from gpu_ppm_amd.lib import Context, PpmError
from oracle_binding import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(data_root, small_nsx_cfg):
    c = Context(data_root, lensing="wide", nsx_path=fixtures.nsx_path(data_root, small=True),
                responses=("std", "std128"), nsx_config=small_nsx_cfg)
    yield c
    c.close()


def cal_chi2(data, model):
    """p_devarshi24st.py:43-49, restated; bins with model = data = 0 add 0
    (the limit model -> 0) instead of 0/0."""
    keep = (model != 0) | (data != 0)
    fac = 1e6 / np.sum(model)
    return np.sum((data[keep] * fac - model[keep] * fac) ** 2 / (model[keep] * fac))


def loglike(data, model):
    keep = (model != 0) | (data != 0)
    d, m = data[keep], model[keep]
    return np.sum(np.where(d == 0, 0.0, d * np.log(np.where(d == 0, 1.0, m))) - m - gammaln(d + 1.0))


def grid(method=abi.METHOD_AUTO):
    return abi.make_grid(64, 16, sampling=abi.SAMPLE_GL, n_gl=2, response=1, method=method)


@pytest.mark.parametrize("method", [abi.METHOD_BATCHED, abi.METHOD_PLAIN])
def test_fit_matches_formulas_on_returned_counts(ctx, method):
    stars = S.sample_priors(5, seed=8)
    g = grid(method)
    print("case: fit, method", method, flush=True)
    _, ref_counts, _ = ctx.eval(stars, g)
    rng = np.random.default_rng(1)
    # one data set per star (the cubic interpolant can undershoot below zero
    # in the Wien tail, as the reference's does: draw from the clipped model)
    data = rng.poisson(np.maximum(ref_counts, 0.0)).astype(np.float64)
    chi2, ll, tot, counts, _ = ctx.eval_fit(stars, g, data, counts=True)
    np.testing.assert_array_equal(counts, ref_counts)          # the same fold
    with np.errstate(invalid="ignore", divide="ignore"):
        for k in range(len(stars)):
            assert tot[k] == pytest.approx(counts[k].sum(), rel=1e-12)
            np.testing.assert_allclose(chi2[k], cal_chi2(data[k], counts[k]), rtol=1e-10, equal_nan=True)
            np.testing.assert_allclose(ll[k], loglike(data[k], counts[k]), rtol=1e-10, equal_nan=True)
    assert np.isfinite(chi2).all()
    # one data array shared by every star = the same array repeated
    shared = data[0]
    c_s, l_s, _, _, _ = ctx.eval_fit(stars, g, shared)
    c_r, l_r, _, _, _ = ctx.eval_fit(stars, g, np.repeat(shared[None], len(stars), axis=0))
    np.testing.assert_array_equal(c_s, c_r)
    np.testing.assert_array_equal(l_s, l_r)


def test_fit_against_oracle_counts(ctx, data_root, small_nsx_cfg):
    st = S.sample_priors(1, seed=9)[0]
    g = grid(abi.METHOD_PLAIN)
    orc = Oracle(data_root, lensing="wide", nsx_path=fixtures.nsx_path(data_root, small=True),
                 responses=("std", "std128"), nsx_cfg=small_nsx_cfg)
    try:
        flux, _ = orc.profile(st, g, threads=16)
        m = orc.fold(flux, 1, st.exposure_s / g.n_phase)
    finally:
        orc.close()
    data = np.random.default_rng(2).poisson(np.maximum(m, 0.0)).astype(np.float64)
    chi2, ll, tot, counts, _ = ctx.eval_fit([st], g, data, counts=True)
    # compare on the bins where both models are positive (the statistics of
    # an undershooting bin are not finite)
    ok = (m > 0) & (counts[0] > 0)
    assert tot[0] == pytest.approx(m.sum(), rel=1e-4)
    assert cal_chi2(data[ok], counts[0][ok]) == pytest.approx(cal_chi2(data[ok], m[ok]), rel=1e-3)
    if ok.all():
        assert chi2[0] == pytest.approx(cal_chi2(data, m), rel=1e-3)
        assert ll[0] == pytest.approx(loglike(data, m), rel=1e-4)


def test_fit_needs_a_response(ctx):
    g = abi.make_grid(32, 8, response=-1)
    with pytest.raises(PpmError) as e:
        ctx.eval_fit(S.sample_priors(1, seed=1), g, np.ones((300, 8)))
    assert e.value.status == abi.ERR_ARG

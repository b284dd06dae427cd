"""Oracle outer (initial-topography) iteration: update rules against a model whose answer is
known, and the warm start against the cold start (run_outer.py:174-279, solver_opt.py:272-286)."""
import numpy as np
import pytest

from oracle.outer import aw_rms, outer_iteration


class _SHTStub:
    def __init__(self, nlat):
        self.weights = np.linspace(0.5, 1.5, nlat)


class _LinearModel:
    """Forward model with tgrid = a T0 + b per cell (a, b fixed): the secant step of two passes
    solves a T0 + b = obs exactly wherever a lies inside [smin, smax]."""

    a = b = None

    def __init__(self, sht, love, T0, ice0, **kw):
        self.T0 = np.array(T0)

    def step(self, ice, dt=None):
        return {"tgrid": self.a * self.T0 + self.b, "numiter": 1}


def test_aw_rms_closed_form():
    w = np.array([1.0, 3.0])
    f = np.array([[2.0, 2.0], [-2.0, -2.0]])
    assert aw_rms(f, w) == pytest.approx(2.0, abs=1e-15)
    g = np.array([[1.0, 1.0], [0.0, 0.0]])
    assert aw_rms(g, w) == pytest.approx(np.sqrt(1.0 / 4.0), abs=1e-15)


def test_picard_and_secant_on_a_linear_model():
    rng = np.random.default_rng(3)
    shape = (6, 12)
    _LinearModel.a = rng.uniform(0.7, 1.3, shape)
    _LinearModel.b = rng.normal(0.0, 50.0, shape)
    obs = rng.normal(-1000.0, 2000.0, shape)
    ice = [np.zeros(shape)] * 3
    kw = dict(lmax=5, dt=1.0, solver_cls=_LinearModel)
    # Picard: T0_{k+1} = T0_k - (a T0_k + b - obs), from T0_1 = obs
    recs, T0, preds = outer_iteration(_SHTStub(6), None, obs, ice, iters=3, outer="picard", **kw)
    t = obs.copy()
    for k in range(3):
        r = _LinearModel.a * t + _LinearModel.b - obs
        assert np.allclose(preds[k] - obs, r, rtol=0, atol=1e-9)
        t = t - r
    assert np.allclose(T0, t, rtol=0, atol=1e-9)
    # secant from pass 2 on: the third pass sits on the fixed point a T0 + b = obs
    recs, T0, preds = outer_iteration(_SHTStub(6), None, obs, ice, iters=3, outer="secant", **kw)
    fixed = (obs - _LinearModel.b) / _LinearModel.a
    assert np.abs(preds[2] - obs).max() < 1e-9 * np.abs(obs).max()
    assert np.allclose(T0, fixed, rtol=1e-12, atol=1e-9)
    assert recs[1]["secant_cells_clamped"] == 0 and recs[1]["secant_cells_picard"] == 0
    # slopes outside [smin, smax] are clamped: with a = 2 everywhere the step is T0 - r / 1.5
    _LinearModel.a = np.full(shape, 2.0)
    recs, T0b, preds = outer_iteration(_SHTStub(6), None, obs, ice, iters=2, outer="secant", **kw)
    assert recs[1]["secant_cells_clamped"] == obs.size
    T0_2 = obs - (2.0 * obs + _LinearModel.b - obs)
    assert np.allclose(T0b, T0_2 - (2.0 * T0_2 + _LinearModel.b - obs) / 1.5, rtol=1e-13, atol=1e-9)


@pytest.mark.slow
def test_warm_start_keeps_the_fixed_point():
    from oracle.love import LoveNumbers
    from oracle.sht import OracleSHT
    from slevolve_amd import synthetic as syn

    w = syn.workload("lmax63")
    sht = OracleSHT(w.nlat, w.nlon, w.lmax)
    love = LoveNumbers(w.love_text, w.lmax, text=True)
    ice = [w.ice(k) for k in range(11)]
    kw = dict(lmax=w.lmax, dt=w.dt, iters=3, epsilon1=1e-10, inner="extrap")
    cold, T0c, pc = outer_iteration(sht, love, w.topo, ice, reuse="off", **kw)
    warm, T0w, pw = outer_iteration(sht, love, w.topo, ice, reuse="warm", **kw)
    # the same fixed point up to the stopping test (a relative change of the L1 norm) and the
    # three-per-step cap on grounded-ice re-guesses, which make the converged state depend
    # slightly on the start: 1.4e-5 m at most here, against misfits of metres
    for a, b in zip(pc, pw):
        assert np.abs(a - b).max() < 1e-4
    assert np.abs(T0c - T0w).max() < 1e-4
    assert warm[2]["numiter"] < cold[2]["numiter"]
    assert cold[1]["rms_pred_minus_obs"] < 0.05 * cold[0]["rms_pred_minus_obs"]

"""GPU outer (initial-topography) iteration against the oracle's (run_outer.py:174-279), with and
without the warm starts across passes (solver_opt.py:272-286,515-528: warm, warm+corr, warm+corr2,
warm+ice), through the C-ABI."""
import numpy as np
import pytest

from oracle.love import LoveNumbers as OLove
from oracle.outer import outer_iteration as oracle_outer
from oracle.sht import OracleSHT
from slevolve_amd import synthetic as syn

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("outer,reuse", [("picard", "off"), ("secant", "warm"), ("picard", "warm+corr"),
                                         ("secant", "warm+corr2"), ("picard", "warm+ice")])
def test_outer_iteration_matches_oracle(outer, reuse):
    from slevolve_amd import SHT, LoveNumbers
    from slevolve_amd.outer import outer_iteration

    w = syn.workload("lmax63")
    ice = [w.ice(k) for k in range(11)]
    kw = dict(lmax=w.lmax, dt=w.dt, iters=3, outer=outer, reuse=reuse, epsilon1=1e-8, inner="extrap")
    print(f"[outer] {outer} {reuse}", flush=True)
    ro, T0o, po = oracle_outer(OracleSHT(w.nlat, w.nlon, w.lmax), OLove(w.love_text, w.lmax, text=True),
                               w.topo, ice, **kw)
    rg, T0g, pg = outer_iteration(SHT(w.nlat, w.nlon, w.lmax), LoveNumbers(w.love_text, w.lmax, text=True),
                                  w.topo, ice, **kw)
    for k, (a, b) in enumerate(zip(rg, ro)):
        assert a["numiter_steps"] == b["numiter_steps"], (k, a["numiter_steps"], b["numiter_steps"])
        err = float(np.abs(pg[k].cpu().numpy() - po[k]).max())
        print(f"[outer] pass {k + 1}: max |pred gpu - oracle| {err:.2e} m, rms {a['rms_pred_minus_obs']:.3e}", flush=True)
        assert err < 1e-8
        for key in ("rms_pred_minus_obs", "max_abs", "rms_T0_minus_obs"):
            assert abs(a[key] - b[key]) <= 1e-8 + 1e-10 * abs(b[key]), (k, key, a[key], b[key])
        for key in ("secant_cells_picard", "secant_cells_clamped"):
            assert a.get(key) == b.get(key)
    assert np.abs(T0g.cpu().numpy() - T0o).max() < 1e-8

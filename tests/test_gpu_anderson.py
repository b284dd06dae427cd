"""GPU parity of the Anderson inner iteration (inner="anderson", slx_step_iterate) against the oracle
(oracle/solver.py + oracle/anderson.py, restating solver_opt.py:67-145,705-775)."""
import numpy as np
import pytest

from oracle.love import LoveNumbers as OLove
from oracle.sht import OracleSHT
from oracle.solver import SeaLevelSolver as OSolver
from slevolve_amd import synthetic as syn

from test_gpu_solver import compare

pytestmark = pytest.mark.gpu


def pair(nlat, L, text, topo, ice0, **kw):
    from slevolve_amd import SHT, LoveNumbers, OptSeaLevelSolver

    a = OptSeaLevelSolver(SHT(nlat, 2 * nlat, L), LoveNumbers(text, L, text=True), topo, ice0, norder=L, **kw)
    b = OSolver(OracleSHT(nlat, 2 * nlat, L), OLove(text, L, text=True), topo, ice0, norder=L, **kw)
    return a, b


def test_small_case_anderson(prem_text):
    from slevolve_amd import SHT

    topo, ice = syn.small_case(SHT(128, 256, 63).gaulats, 256)
    a, b = pair(128, 63, prem_text, topo, ice[0], dt=1000.0, epsilon1=1e-5, inner="anderson")
    for k in range(1, 9):
        print(f"[anderson small case] step {k}", flush=True)
        compare(a.step(ice[k]), b.step(ice[k]))
    assert a.inner_stats == b.inner_stats, (a.inner_stats, b.inner_stats)
    assert a.gr_counts == b.gr_counts


def test_lmax63_workload_anderson_counts():
    w = syn.workload("lmax63")
    a, b = pair(w.nlat, w.lmax, w.love_text, w.topo, w.ice(0), dt=w.dt, epsilon1=1e-5, inner="anderson")
    na, nb = [], []
    for k in range(1, 31):
        ra, rb = a.step(w.ice(k)), b.step(w.ice(k))
        na.append(ra["numiter"])
        nb.append(rb["numiter"])
        assert float(np.abs(ra["tgrid"].cpu().numpy() - rb["tgrid"]).max()) < 1e-9, k
    print("numiter", na, a.inner_stats, flush=True)
    assert na == nb and sum(na) == 176
    assert a.inner_stats == b.inner_stats


@pytest.mark.parametrize("depth,freeze,keep", [(1, 2, False), (5, 3, True)])
def test_anderson_options_against_oracle(prem_text, topo_seed, depth, freeze, keep):
    lat, lon, topo, ice = syn.synthetic_earth(64, topo_seed)
    kw = dict(dt=500.0, tpw=True, checkmarine=True, calcRG=True, epsilon1=1e-7, inner="anderson",
              anderson_depth=depth, anderson_freeze=freeze, anderson_keep=keep, max_mask_refresh=1)
    a, b = pair(64, 63, prem_text, topo, ice[0], **kw)
    for k in range(1, 7):
        print(f"[anderson options {depth},{freeze},{keep}] step {k}", flush=True)
        compare(a.step(ice[k]), b.step(ice[k]))
    assert a.inner_stats == b.inner_stats

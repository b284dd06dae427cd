"""GPU parity of the second-order load placement (ramp), the analytic rates and the horizontal
read-out (slx_step_finish_readouts) against the oracle, and their invariants at full size.

Tolerances: topography, R, G 1e-9 m as for the default step; rate grids and spectra 1e-11 m/yr
(the rates are ~0.1 m/yr, so this is ~1e-10 relative); rotational rates 1e-9 relative; the
horizontal spectrum 1e-10 relative to its largest coefficient."""
import numpy as np
import pytest

from oracle.love import LoveNumbers as OLove
from oracle.sht import OracleSHT
from oracle.solver import Planet
from oracle.solver import SeaLevelSolver as OSolver
from slevolve_amd import synthetic as syn

from test_gpu_solver import compare

pytestmark = pytest.mark.gpu

VSTEPS = [1000.0, 500.0, 2000.0, 250.0, 1000.0, 750.0, 1500.0, 500.0]


def pair(nlat, L, text, topo, ice0, horizontal=False, **kw):
    from slevolve_amd import SHT, LoveNumbers, OptSeaLevelSolver

    a = OptSeaLevelSolver(SHT(nlat, 2 * nlat, L), LoveNumbers(text, L, text=True), topo, ice0, norder=L, **kw)
    b = OSolver(OracleSHT(nlat, 2 * nlat, L), OLove(text, L, text=True), topo, ice0, norder=L, **kw)
    a.horizontal = b.horizontal = horizontal
    return a, b


def compare_readouts(ra, rb, rates=True, horizontal=False):
    h = lambda t: t.cpu().numpy() if hasattr(t, "cpu") else np.asarray(t)  # noqa: E731
    if rates:
        x, y = ra["rates"], rb["rates"]
        for k in ("dSLdt", "dRdt", "dSLdt_lm", "dRdt_lm"):
            e = float(np.abs(h(x[k]) - y[k]).max())
            assert e < 1e-11, (k, e)
        if "mdot" in y:
            for k in ("ildot", "mdot", "lambdadot"):
                scale = float(np.abs(y[k]).max())
                assert float(np.abs(x[k] - y[k]).max()) <= 1e-9 * scale + 1e-300, (k, x[k], y[k])
    if horizontal:
        d, o = h(ra["horiz"]), rb["horiz"]
        assert float(np.abs(d - o).max()) < 1e-10 * float(np.abs(o).max())


@pytest.mark.parametrize("ramp", [True, False])
def test_quickstart_case_readouts(prem_text, topo_seed, ramp):
    """Non-axisymmetric caps: every rotational term (orders 0-2) is exercised."""
    lat, lon, topo, ice = syn.synthetic_earth(128, topo_seed)
    a, b = pair(128, 63, prem_text, topo, ice[0], horizontal=True, dt=1000.0, tpw=True, checkmarine=True,
                calcRG=True, epsilon1=1e-5, ramp=ramp)
    for k in range(1, 9):
        print(f"[quickstart readouts ramp={ramp}] step {k}", flush=True)
        ra, rb = a.step(ice[k], rates=True), b.step(ice[k], rates=True)
        compare(ra, rb)
        compare_readouts(ra, rb, horizontal=True)
    assert abs(rb["rates"]["mdot"][0]) > 0


def test_small_case_ramp_variable_steps(prem_text):
    from slevolve_amd import SHT

    topo, ice = syn.small_case(SHT(128, 256, 63).gaulats, 256)
    a, b = pair(128, 63, prem_text, topo, ice[0], dt=VSTEPS[0], epsilon1=1e-5, ramp=True)
    for k in range(1, 9):
        print(f"[small case ramp vardt] step {k}", flush=True)
        ra, rb = a.step(ice[k], dt=VSTEPS[k - 1], rates=True), b.step(ice[k], dt=VSTEPS[k - 1], rates=True)
        compare(ra, rb)
        compare_readouts(ra, rb)


def test_readouts_leave_the_step_unchanged(prem_text, topo_seed):
    """Rates and the horizontal read-out are output only: the step is bit-identical without them."""
    from slevolve_amd import SHT, LoveNumbers, OptSeaLevelSolver

    lat, lon, topo, ice = syn.synthetic_earth(64, topo_seed)
    kw = dict(norder=63, dt=500.0, epsilon1=1e-7, ramp=True)
    a = OptSeaLevelSolver(SHT(64, 128, 63), LoveNumbers(prem_text, 63, text=True), topo, ice[0], **kw)
    b = OptSeaLevelSolver(SHT(64, 128, 63), LoveNumbers(prem_text, 63, text=True), topo, ice[0], **kw)
    b.horizontal = True
    for k in range(1, 6):
        ra, rb = a.step(ice[k]), b.step(ice[k], rates=True)
        for key in ("tgrid", "R", "G", "dS_converged"):
            assert np.array_equal(ra[key].cpu().numpy(), rb[key].cpu().numpy()), key
        assert ra["numiter"] == rb["numiter"]
        assert ra["rates"] is None and ra["horiz"] is None and rb["rates"] is not None


def test_lmax512_ramp_rates_against_oracle():
    """The 512 x 1024, lmax 512 workload (30 modes per degree, marine sector) with the ramp, rates
    and the horizontal read-out: 3 steps against the oracle, and the degree-0 ice balance of the
    rate grid on the device: analysis(C* dSL/dt)_00 = -(rho_i/rho_w) dI*_00 / dt up to the legacy
    degree-0 normalisation."""
    from slevolve_amd import SHT, LoveNumbers, OptSeaLevelSolver

    w = syn.workload("lmax512")
    kw = dict(norder=w.lmax, dt=w.dt, tpw=True, checkmarine=True, calcRG=True, epsilon1=1e-5, ramp=True)
    sht = SHT(w.nlat, w.nlon, w.lmax)
    a = OptSeaLevelSolver(sht, LoveNumbers(w.love_text, w.lmax, text=True), w.topo, w.ice(0), **kw)
    b = OSolver(OracleSHT(w.nlat, w.nlon, w.lmax), OLove(w.love_text, w.lmax, text=True), w.topo, w.ice(0), **kw)
    a.horizontal = b.horizontal = True
    p = Planet()
    for k in range(1, 4):
        print(f"[lmax512 ramp readouts] step {k}", flush=True)
        before = a.state.prev_ice[0, 0].item()
        ra, rb = a.step(w.ice(k), rates=True), b.step(w.ice(k), rates=True)
        compare(ra, rb)
        compare_readouts(ra, rb, horizontal=True)
        if k >= 2:
            dice00 = (a.state.prev_ice[0, 0].item() - before).real
            rhs = -(p.rhoi / p.rhow) * dice00 / w.dt
            cst = a.cxy * a.beta
            lhs = sht.analysis(ra["rates"]["dSLdt"] * cst)[0, 0].real.item()
            c00 = sht.analysis(cst)[0, 0].real.item()
            d00 = ra["rates"]["dSLdt_lm"][0, 0].real.item()
            assert abs(lhs - rhs) <= 2e-9 * abs(d00 * c00) + 1e-14 * abs(rhs), (lhs, rhs)
